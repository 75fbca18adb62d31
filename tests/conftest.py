import json
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

H2O = 'O 0 0 0; H 0 -0.757 0.587; H 0 0.757 0.587'


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: needs a HIP device (MI355X)')
    config.addinivalue_line('markers', 'slow: longer CPU test')


@pytest.fixture(scope='session')
def pins():
    with open(os.path.join(ROOT, 'tests', 'golden', 'reference_pins.json')) as f:
        return json.load(f)


@pytest.fixture(scope='session')
def oracle():
    from oracle import oracle as O
    O.build()
    return O


@pytest.fixture(scope='session')
def h2o_vdz():
    from pyscf_amd import gto
    return gto.M(atom=H2O, basis='cc-pvdz')


@pytest.fixture(scope='session')
def h2o_631g():
    from pyscf_amd import gto
    return gto.M(atom=H2O, basis='631g')
