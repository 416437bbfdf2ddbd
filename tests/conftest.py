import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a gfx950 GPU (run with -m gpu)")


def golden_bytes(name: str) -> bytes:
    with open(os.path.join(GOLDEN, name), "rb") as f:
        return f.read()


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def hip_ctx():
    """A HIP context for GPU tests. A GPU test never skips: on a GPU run a
    missing device is a failure."""
    import fritillaria_amd as fa
    try:
        return fa.HipContext(0)
    except Exception as exc:  # pragma: no cover - only on a broken GPU box
        pytest.fail(f"GPU test without a usable device: {exc}")


@pytest.fixture(scope="session")
def codec(hip_ctx):
    import fritillaria_amd as fa
    return fa.HipCodec(hip_ctx)


@pytest.fixture(scope="session")
def decoder(hip_ctx):
    import fritillaria_amd as fa
    return fa.BamDecoder(hip_ctx)
