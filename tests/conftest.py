import json
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_ROOT = os.path.join(ROOT, "dreams-rns-cuda_amd")
for p in (PKG_ROOT, ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden", "reference_vectors.json")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libdreams_hip.so")


@pytest.fixture(scope="session")
def golden():
    with open(GOLDEN) as f:
        return json.load(f)


def program_from_json(d):
    from dreams_rns.compiler import CmfProgram, Instruction, Opcode

    return CmfProgram(d["m"], d["dim"], [Instruction(Opcode(o[0]), o[1], o[2], o[3]) for o in d["ops"]],
                      [int(c) for c in d["constants"]], list(d["directions"]))


@pytest.fixture(scope="session")
def gpu_ctx():
    from dreams_rns import _native as N

    try:
        return N.context()
    except N.NativeUnavailable as e:
        pytest.fail(f"GPU test without the native engine: {e}")
