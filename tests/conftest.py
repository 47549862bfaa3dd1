import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) to run")


def has_gpu():
    try:
        import emptyheaded_amd as E
        return E.device_count() > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.lib()
    return O


@pytest.fixture(scope="session")
def ctx():
    import emptyheaded_amd as E
    if E.device_count() == 0:
        pytest.fail("no GPU visible: -m gpu tests need an MI355X")
    c = E.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="session")
def facebook():
    import numpy as np
    d = np.load(os.path.join(ROOT, "tests", "golden", "facebook_pruned.npz"))
    return d["src"], d["dst"]


@pytest.fixture(scope="session")
def facebook_dup():
    import numpy as np
    d = np.load(os.path.join(ROOT, "tests", "golden", "facebook_duplicated.npz"))
    return d["src"], d["dst"]


@pytest.fixture(scope="session")
def fixtures_meta():
    import json
    with open(os.path.join(ROOT, "tests", "golden", "fixtures.json")) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def synthetic_meta():
    import json
    with open(os.path.join(ROOT, "tests", "golden", "synthetic.json")) as f:
        return json.load(f)
