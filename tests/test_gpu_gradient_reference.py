"""The energy's autograd Function (its hand-written backward: the adjoint
2 H psi dv, the convolution's adjoint reusing the symmetric kernel spectrum)
against the reference's own torch autograd gradient of its energy
(tests/golden/gradient_reference.npz) on the six physics sets, to 1e-12
relative."""
import pytest

from tests.golden.make_operator_golden import CASES, state
from tests.helpers import cfg
from tests.test_oracle_gradient_reference import reference_gradient
from tests.test_oracle_operators_reference import close_vec

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.mark.parametrize("i", range(len(CASES)))
def test_energy_backward_matches_reference_autograd(i):
    import becgpp_amd as B
    c = cfg(**CASES[i])
    G = B.make_grid(c)
    print(f"[gradient] reference autograd case {i}: {CASES[i]}", flush=True)
    p = torch.as_tensor(state(CASES[i], seed=700 + i)).cuda().requires_grad_(True)
    B.energy_t(p, G, c["beta2"], c["beta3"], c["Omega"], c["G_C"]).backward()
    close_vec(p.grad.detach().cpu().numpy(), reference_gradient(i))
