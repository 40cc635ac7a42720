import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def load_golden(name):
    return dict(np.load(os.path.join(GOLDEN, name), allow_pickle=False))


def mesh_from_golden(d):
    from fast_amd.mesh import Mesh
    kw = dict(coords=d["coords"], conn=d["conn"], shape=str(d["shape"]),
              nip=tuple(int(x) for x in d["nip"]), form=str(d["form"]),
              mat=d["mat"], fixed=d["fixed"], deg=tuple(int(x) for x in d["deg"]))
    if "ext" in d:
        kw.update(ext=d["ext"], cpw=d["cpw"], span=d["span"])
    if "fext" in d:
        kw.update(fext=d["fext"])
    return Mesh(**kw)


@pytest.fixture(scope="session")
def oracle_mod():
    from oracle import oracle
    oracle.lib()
    return oracle
