import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) to run")


def pytest_collection_modifyitems(config, items):
    # GPU tests run wherever a ROCm device node exists; on a host without one they are skipped
    # rather than failing on device allocation.
    if os.path.exists("/dev/kfd"):
        return
    skip = pytest.mark.skip(reason="no ROCm device on this host")
    for item in items:
        if "gpu" in item.keywords:
            item.add_marker(skip)


@pytest.fixture(scope="session")
def known_answers():
    with open(os.path.join(GOLDEN, "known_answers.json")) as fh:
        return json.load(fh)


@pytest.fixture(scope="session")
def snapshots():
    return dict(np.load(os.path.join(GOLDEN, "reference_snapshots.npz")))


def golden_system(name, **kw):
    """(topology, coordinates) of a reference test system: 'trpcage' or 'ala_dipeptide'."""
    from stormm_amd.topology import read_inpcrd, read_prmtop
    top = os.path.join(GOLDEN, "inputs", f"{name}.top")
    crd = os.path.join(GOLDEN, "inputs", f"{name}.inpcrd")
    return read_prmtop(top, **kw), read_inpcrd(crd)
