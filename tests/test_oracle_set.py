"""The reader of the reference's oracle dump sets (`mccuda-oracle/1` + `mccuda-terrain-oracle/1`)
on a set written from the oracle in that layout: shapes, byte orders, hashes, errors; the
named-function dumps against the oracle's volume sampler (the reference's L1 check, run on the
CPU as the reference runs it)."""
import json
import os

import numpy as np
import pytest

import oracle_data
from oracle_set import write_set

SEEDS = [(123456789, [(0, 0), (-3, 7)]), (-1, [(6128, 0)])]
NAMED = ['final_density', 'climate/temperature']


@pytest.fixture(scope='module')
def oset(orc, tmp_path_factory):
    d = str(tmp_path_factory.mktemp('oracle_set'))
    write_set(d, orc, SEEDS, named=NAMED)
    return d


def test_reader_shapes_and_hashes(oset, orc):
    s = oracle_data.load_oracle_set(oset)
    assert (s.min_y, s.height) == (-64, 384) and len(s.seeds) == 2
    assert [c.tag for c in s.seeds[0].chunks] == ['c0', 'c1'] and s.seeds[1].seed == -1
    assert s.verify_hashes() == []
    d, f = s.density(s.seeds[0]), s.fill(s.seeds[0])
    b, a, t = s.terrain(s.seeds[0])
    assert d.shape == (2, 384, 16, 16) and d.dtype == np.float32 and f.shape == (2, 384, 16, 16)
    assert b.shape == (2, 96, 4, 4) and a.dtype == np.uint16 and t.shape == (2, 384, 16, 16)
    ot, ob, oa, od = orc.generate(123456789, [(0, 0), (-3, 7)], density=True)
    assert (d.view(np.uint32) == od.view(np.uint32)).all() and (t == ot).all() and (a == oa).all() and (b == ob).all()


def test_named_function_dumps_match_the_oracle_sampler(oset, orc):
    s = oracle_data.load_oracle_set(oset)
    ir = json.load(open(__import__('oracle').IR_PATH))
    n = 0
    for seed in s.seeds:
        for dump in seed.node_dumps:
            k = int(dump.split('_c')[1])
            c = seed.chunks[k]
            for name, file in s.named:
                want = s.node(dump, file)
                got = orc.sample(seed.seed, ir['roots'][name], [16, 384, 16, c.cx * 16, -64, c.cz * 16, 1, 1, 1]).reshape(16, 16, 384).transpose(2, 0, 1)
                assert (want.view(np.uint32) == np.ascontiguousarray(got).view(np.uint32)).all(), (dump, name)
                n += 1
    assert n == 4


def test_reader_errors(oset, tmp_path):
    s = oracle_data.load_oracle_set(oset)
    with pytest.raises(oracle_data.OracleError):
        oracle_data.load_oracle_set(str(tmp_path))                       # no manifest
    m = json.load(open(os.path.join(oset, 'manifest.json')))
    m['format'] = 'mccuda-oracle/2'
    json.dump(m, open(tmp_path / 'manifest.json', 'w'))
    with pytest.raises(oracle_data.OracleError):
        oracle_data.load_oracle_set(str(tmp_path))                       # unknown format
    bad = tmp_path / 'copy'
    import shutil
    shutil.copytree(oset, bad)
    with open(bad / 'fill' / 's1.u8', 'r+b') as fh:                      # one changed byte
        fh.seek(1000)
        v = fh.read(1)
        fh.seek(1000)
        fh.write(bytes([v[0] ^ 1]))
    assert oracle_data.load_oracle_set(str(bad)).verify_hashes() == ['fill/s1.u8']
    with open(bad / 'density' / 's0.f32', 'ab') as fh:                   # wrong size
        fh.write(b'\0\0\0\0')
    with pytest.raises(oracle_data.OracleError):
        oracle_data.load_oracle_set(str(bad)).density(s.seeds[0])
