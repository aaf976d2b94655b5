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


def ref_lib():
    """The reference cpu/ port compiled from its sources (oracle/_ref), or None."""
    path = os.path.join(ROOT, 'oracle', '_ref', 'libmcref.so')
    if not os.path.exists(path):
        return None
    P, I64, I = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int
    R = ctypes.CDLL(path)
    R.ref_create.restype = P
    R.ref_create.argtypes = [ctypes.c_char_p, ctypes.c_char_p]
    R.ref_last_error.restype = ctypes.c_char_p
    R.ref_generate_terrain.argtypes = [P, I64, P, I, P, P, P, P]
    R.ref_generate_region.argtypes = [P, I64, I, I, I, I, I, P, P]
    R.ref_seed_layers.argtypes = [P, I64, P, P]
    R.ref_noise_ref.argtypes = [P, I64, ctypes.c_char_p, P, P]
    R.ref_seed_misc.argtypes = [P, I64, P, P, P]
    R.ref_tree_search.argtypes = [P, P, P, I, P]
    R.ref_zoom.argtypes = [I64, P, I, P]
    R.ref_sample.argtypes = [P, I64, ctypes.c_char_p, P, P]
    R.ref_point.argtypes = [P, I64, ctypes.c_char_p, P, I, P]
    return R


@pytest.fixture(scope='session')
def ref():
    R = ref_lib()
    if R is None:
        pytest.skip('reference cpu/ port not built (oracle/_ref)')
    import mct
    h = R.ref_create(mct.IR_PATH.encode(), mct.TERRAIN_PATH.encode())
    assert h, R.ref_last_error()
    return R, ctypes.c_void_p(h)
