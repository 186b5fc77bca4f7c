"""The HIP operators against the reference's own applications
(tests/golden/operator_reference.npz; the oracle meets the same vectors in
tests/test_oracle_operators_reference.py): H psi, every energy component with
phi, the residual triple, L_z psi and the kernel convolution for the six
physics sets of the reference's FD audit, each to 1e-12 relative."""
import pytest

from tests.golden.make_operator_golden import CASES, state
from tests.helpers import cfg
from tests.test_oracle_operators_reference import check_operators

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.mark.parametrize("i", range(len(CASES)))
def test_hip_operators_match_reference(i):
    import becgpp_amd as B
    c = cfg(**CASES[i])
    G = B.make_grid(c)
    p = torch.as_tensor(state(CASES[i], seed=700 + i)).cuda()
    print(f"[operators] reference case {i}: {CASES[i]}", flush=True)
    comp = B.energy_components(p, G, c["beta2"], c["beta3"], c["Omega"], c["G_C"])
    host = {k: float(v) for k, v in comp.items() if k != "phi"}
    phi = comp["phi"].cpu().numpy() if torch.is_tensor(comp["phi"]) else None
    conv = B.kernel_convolution(p.abs() ** 2, G).cpu().numpy() if G["kernel"] != "none" else None
    check_operators(i, B.apply_H(p, G, c).cpu().numpy(), host, phi, B.compute_residual(p, G, c),
                    B.Lz_op(p, G).cpu().numpy(), conv)
