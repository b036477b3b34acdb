"""The oracle's fd::gradient (render + MseToTarget objective) pinned by the
adjoint gradient and the reference's FD semantics, and its Adam epoch
pinned by the reference's Adam tests and the first-epoch closed form."""
import pytest

import fd_cases as fc


@pytest.mark.parametrize("case", fc.FD_CASES, ids=lambda f: f.__name__)
def test_fd_case_oracle(orc, case):
    case(orc.fd_gradient, orc)


@pytest.mark.parametrize("case", fc.ADAM_CASES, ids=lambda f: f.__name__)
def test_adam_case_oracle(orc, case):
    case(orc.adam_step)
