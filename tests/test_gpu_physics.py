"""The reference's physics/convention cases (physics_cases.py) on the HIP
engines: the f64 host drop-in and complex128 device form at the f64 bars, the
FP32-bulk complex64 form with tolerances floored at C64_FLOOR."""
import pytest

import physics_cases as pc

pytestmark = pytest.mark.gpu

ALL = pc.ASM_CASES + pc.FRESNEL_CASES + pc.RS_CASES + pc.FRAUNHOFER_CASES + pc.MASK_CASES
C64_FLOOR = 2e-6  # about 16 FP32 ulps on unit-scale quantities


@pytest.fixture(scope="module")
def engines():
    import optics_gpu as G
    return {"f64": (G, 1.0), "c128": (G.DevEngine("c128"), 1.0), "c64": (G.DevEngine("c64"), C64_FLOOR)}


@pytest.mark.parametrize("kind", ["f64", "c128", "c64"])
@pytest.mark.parametrize("case", ALL, ids=[c.__name__ for c in ALL])
def test_gpu_case(engines, kind, case):
    eng, tol = engines[kind]
    print(kind, case.__name__, flush=True)
    case(eng, tol)
