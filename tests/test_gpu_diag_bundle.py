"""The device diagnostics (becgpp_amd.diagnostics: observables, radii, vortex
counts, LLL weight, TF fields) and resampling against the bundles and states
the reference computed itself (tests/golden/diag_reference.npz; the oracle
meets the same vectors in tests/test_oracle_diag_bundle.py).  Radii picked by
a cumulative sum or a bin may sit one grid step apart where the device's sums,
added in another order, tie at rounding size; every other value to 1e-11 (TF
fields 1e-9, resampled states 1e-13)."""
import numpy as np
import pytest

from tests.golden.make_diag_golden import CASES, RESAMPLE, state
from tests.helpers import cfg
from tests.test_oracle_diag_bundle import GOLD, check_bundle

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def device_state(i, G):
    kw, zeros = CASES[i]
    psi = torch.as_tensor(state(kw, zeros, seed=500 + i))
    return (psi / torch.sqrt((psi.abs() ** 2).sum() * float(G["dx"]) ** int(G["ndim"]))).cuda()


@pytest.mark.parametrize("i", range(len(CASES)))
def test_device_bundle_matches_reference(i):
    import becgpp_amd as B
    c = cfg(tf_cache=False, **CASES[i][0])
    G = B.make_grid(c)
    print(f"[diag] reference bundle case {i}: {CASES[i][0]}", flush=True)
    check_bundle(B.diagnostics(device_state(i, G), G, c), i, float(G["dx"]), grid_step_radii=True)


@pytest.mark.parametrize("j", range(len(RESAMPLE)))
def test_device_resample_matches_reference(j):
    import becgpp_amd as B
    i, L2, N2 = RESAMPLE[j]
    c = cfg(**CASES[i][0])
    G = B.make_grid(c)
    G2 = B.make_grid(dict(c, L=L2, Ngrid=N2))
    print(f"[diag] reference resample {j}: case {i} to L={L2} N={N2}", flush=True)
    got = B.resample_state(device_state(i, G), G, G2).cpu().numpy()
    ref = np.load(GOLD)[f"resampled{j}"]
    assert np.abs(got - ref).max() <= 1e-13 * np.abs(ref).max()
