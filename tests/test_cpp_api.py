"""The C++ face of the library (include/mct_terrain_gpu.hpp): it compiles and links against
libmct.so (CPU), and on an MI355X a C++ caller gets the oracle's bytes (GPU)."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from conftest import PKG, ROOT

SRC = os.path.join(ROOT, 'tests', 'cpp', 'generator_example.cpp')


def build(tmp):
    cxx = shutil.which('g++')
    if cxx is None or not os.path.exists(os.path.join(PKG, 'libmct.so')):
        pytest.skip('g++ or libmct.so missing')
    exe = os.path.join(tmp, 'generator_example')
    subprocess.run([cxx, '-std=c++17', '-O1', '-I', os.path.join(ROOT, 'include'), SRC, '-o', exe, '-L', PKG, '-lmct', '-Wl,-rpath,' + PKG],
                   check=True)
    return exe


def test_cpp_class_compiles_and_links(tmp_path):
    assert os.path.exists(build(str(tmp_path)))


@pytest.mark.gpu
def test_cpp_caller_gets_oracle_bytes(tmp_path, orc):
    import mct
    exe = build(str(tmp_path))
    out = os.path.join(str(tmp_path), 'out.bin')
    chunks = [(0, 0), (-1, 5), (6128, 0)]
    args = [exe, mct.IR_PATH, mct.TERRAIN_PATH, '-1', out] + [str(v) for c in chunks for v in c]
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    data = np.fromfile(out, np.uint8)
    n = len(chunks)
    t, b = data[:n * 98304].reshape(n, 384, 16, 16), data[n * 98304:].reshape(n, 96, 4, 4)
    ot, ob, _ = orc.generate(-1, chunks)
    assert np.array_equal(t, ot) and np.array_equal(b, ob)
