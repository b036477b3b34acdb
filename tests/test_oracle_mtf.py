"""The oracle's analysis::mtf pinned by the reference's MTF tests, closed
forms and numpy's fft2."""
import pytest

import mtf_cases as mc


@pytest.mark.parametrize("case", mc.CASES, ids=lambda f: f.__name__)
def test_mtf_case_oracle(orc, case):
    case(orc.mtf)
