import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def golden(name):
    import numpy as np
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def system_from_golden(g):
    import numpy as np
    from ndyn_omp_amd.workloads import System
    ipos = np.array(g["ipos"]) if "ipos" in g.files else None
    init_box = np.array(g["init_box"]) if "init_box" in g.files else None
    return System(pos=np.array(g["pos"]), vel=np.array(g["vel"]), typ=np.array(g["typ"], np.int32),
                  box=np.array(g["box"]), masses=np.array(g["masses"]),
                  pairmat=np.array(g["pairmat"], np.int32), kinds=np.array(g["kinds"], np.int32),
                  params=np.array(g["params"]), ptrmin=float(g["ptrmin"]),
                  ptsize=int(g["ptsize"]), ipos=ipos, init_box=init_box)


@pytest.fixture(scope="session")
def oracle():
    from oracle import pyoracle
    pyoracle.lib()
    return pyoracle
