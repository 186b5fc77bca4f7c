"""The batched multi-seed solve (ground_state_multiseed: the seeds of a point
as one batch) against the reference's own selections with the seeds' noise
switched off (tests/golden/multiseed_reference.json; the oracle meets them in
tests/test_oracle_multiseed.py): every seed's energy, the winner, the spread
and the selected E, mu and L_z."""
import pytest

from tests.golden.make_multiseed_golden import CASES, case_cfg
from tests.helpers import cfg
from tests.test_oracle_multiseed import check_selection

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.mark.parametrize("i", range(len(CASES)))
def test_batched_multiseed_matches_reference(i):
    import becgpp_amd as B
    c = cfg(tf_cache=False, **case_cfg(CASES[i]))
    print(f"[multiseed] reference selection case {i}: {CASES[i]}", flush=True)
    _, _, obs = B.ground_state_multiseed(c, B.make_grid(c))
    check_selection(obs, i)
