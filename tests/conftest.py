"""Test configuration: the `gpu` marker, import paths, and loaders for the checkers.

CPU tests (`-m "not gpu"`) cover the oracle against the committed golden fixtures and
known answers, the code generator's model tables, and that libmct.so exports the C ABI.
GPU tests (`-m gpu`) run the HIP library through the C ABI and compare it bit for bit with
the oracle (and with the reference cpu/ port when its build is present).
"""
import ctypes
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, 'minecraft-terrain-cuda_amd')
GOLDEN = os.path.join(ROOT, 'tests', 'golden')
sys.path.insert(0, PKG)
sys.path.insert(0, os.path.join(ROOT, 'oracle'))


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: needs an MI355X (HIP library calls)')


def _oracle_module():
    import oracle
    return oracle


@pytest.fixture(scope='session')
def orc():
    """The oracle (oracle/liboracle.so) over the synthetic overworld model."""
    return _oracle_module().Oracle()


@pytest.fixture(scope='session')
def golden():
    """Reference outputs written by tests/golden/make_golden.py."""
    import numpy as np
    with np.load(os.path.join(GOLDEN, 'fixtures.npz'), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.fixture(scope='session')
def reference():
    """The reference cpu/ port compiled from its sources (oracle/_ref), where built."""
    m = _oracle_module()
    if not os.path.exists(m.REF_PATH):
        pytest.skip('reference cpu/ port not built (make -C oracle ref)')
    return m.Reference()
