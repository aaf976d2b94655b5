"""The C ABI library loads and exports every function include/mct.h declares (no GPU calls)."""
import ctypes
import os
import re

import pytest

from conftest import PKG, ROOT


def declared_functions():
    text = open(os.path.join(ROOT, 'include', 'mct.h')).read()
    text = re.sub(r'/\*.*?\*/', '', text, flags=re.S)
    return sorted(set(re.findall(r'\b(mct_[a-z_0-9]+)\s*\(', text)))


def test_header_declares_the_boundary():
    names = declared_functions()
    for required in ('mct_create', 'mct_destroy', 'mct_set_seed', 'mct_generate', 'mct_generate_multi', 'mct_biome_tie_stats',
                     'mct_last_batch_done_ms', 'mct_alloc_pinned', 'mct_free_pinned', 'mct_last_error'):
        assert required in names


def test_library_exports_every_declared_symbol():
    path = os.path.join(PKG, 'libmct.so')
    if not os.path.exists(path):
        pytest.skip('libmct.so not built')
    lib = ctypes.CDLL(path)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
