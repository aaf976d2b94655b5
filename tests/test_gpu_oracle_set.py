"""The oracle-set checker (the reference's L2 and terrain checks, tests/test_chunks.cpp and
tests/test_terrain.cpp) run against the HIP library on a set written from the oracle in the
reference's dump layout: every check passes; one changed value in the expected data is
reported as exactly one mismatch (and its file's hash as changed)."""
import os

import numpy as np
import pytest

import oracle_check
from oracle_set import write_set

pytestmark = pytest.mark.gpu

SEEDS = [(123456789, [(0, 0), (1, 0), (-2048, -2048)]), (-1, [(1874999, -1874999), (11, 6128)]), (42, [(-3, -6129)])]


def test_library_passes_the_oracle_set_checks(orc, tmp_path):
    import mct
    mct.lib()
    d = str(tmp_path / 'set')
    write_set(d, orc, SEEDS)
    ok, checks, bad = oracle_check.run(d)
    assert ok and not bad
    assert checks['density'].total == 6 * 98304 and checks['terrain'].total == 6 * 98304 and checks['biomes'].total == 6 * 1536
    # one changed value in the expected terrain and in the expected density
    t = np.fromfile(os.path.join(d, 'terrain', 'terrain_s1.u16'), '<u2')
    t[98304 + 5000] ^= 1
    t.tofile(os.path.join(d, 'terrain', 'terrain_s1.u16'))
    dens = np.fromfile(os.path.join(d, 'density', 's0.f32'), '<f4')
    dens[777] = np.nextafter(dens[777], np.float32(np.inf))
    dens.tofile(os.path.join(d, 'density', 's0.f32'))
    ok2, checks2, bad2 = oracle_check.run(d)
    assert not ok2 and sorted(bad2) == ['density/s0.f32', 'terrain/terrain_s1.u16']
    assert checks2['terrain'].failed == 1 and checks2['density'].failed == 1
    assert checks2['core fill'].failed == 0 and checks2['biomes'].failed == 0 and checks2['aquifer fill'].failed == 0
