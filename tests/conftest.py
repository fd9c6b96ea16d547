import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT / "hyper-vlasov_amd"))
sys.path.insert(0, str(ROOT / "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    # load libhvamd.so (and the ROCm runtime it is built against) before any
    # test module imports torch, so the process holds one HIP runtime
    import hvamd
    try:
        hvamd.load()
    except RuntimeError:
        pass  # not built: the tests that need it fail loudly on their own


@pytest.fixture(scope="session")
def oracle():
    from oracle_lib import Oracle
    return Oracle()


@pytest.fixture(scope="session")
def ref():
    from oracle_lib import Reference
    if not Reference.available:
        pytest.skip("oracle/_ref/libhvref.so not built (needs the reference checkout)")
    return Reference()
