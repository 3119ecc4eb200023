import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

# Worker threads for the oracle: a one-GPU job is granted 16 host cores.
ORACLE_THREADS = int(os.environ.get("RAYCER_ORACLE_THREADS", "16"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a gfx950 GPU (run on the MI355X box)")
    config.addinivalue_line("markers", "slow: long-running (full-size workloads)")


@pytest.fixture(scope="session")
def ctx():
    """One HIP context for the whole GPU session (fails loudly without a GPU)."""
    import raycer_amd as ra

    c = ra.Context(0)
    yield c
    c.close()


@pytest.fixture(params=["fp32", "fp64"])
def decision(request, ctx):
    """Runs a test under each triangle decision mode (RC_HIT_FP32, the
    default, and RC_HIT_FP64); yields True for FP32."""
    ctx.set_hit_decision(request.param)
    yield request.param == "fp32"
    ctx.set_hit_decision("fp32")
