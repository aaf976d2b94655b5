"""The workload definitions the bench, the GPU tests and the fixture writer share."""
import numpy as np

import workloads


def test_fold_chunks_include_the_reference_wrap_points():
    """The reference's noise-wrap chunks (6128, 0) and (-3, -6129) are the folds of the base 3D
    noise's product 684.412 x 0.25 = 171.103 (Corpus.java:30-34): the fold set must hold a chunk
    at x = 6128 and one at z = -6129, and every fold chunk must hold its fold block."""
    prods = workloads.fold_products()
    assert any(abs(p - 171.103) < 1e-9 for p in prods)
    ch = workloads.fold_chunks()
    assert any(x == 6128 for x, _ in ch) and any(z == -6129 for _, z in ch)
    assert len(set(ch)) == len(ch)
    for p in prods:
        b = workloads.FOLD / p
        if b < 16 * 1874999:
            assert any(16 * x <= b < 16 * x + 16 for x, _ in ch)
            assert any(16 * z <= -b < 16 * z + 16 for _, z in ch)


def test_border_region_is_the_reference_verify_region():
    r = workloads.border_region()
    assert len(r) == 64 and min(r) == (1874992, 1874992) and max(r) == (1874999, 1874999)


def test_stage_units_count_distinct_positions():
    """gen_host.stage_units against a brute-force count on two seeds' overlapping regions."""
    import gen_host
    ch = workloads.region(0, 0, 3) + workloads.region(1, 1, 2) + workloads.region(-5, 7, 2)
    sd = np.array([1] * 9 + [2] * 4 + [1] * 4)
    bc, cols, gk = set(), set(), set()
    for s, (x, z) in zip(sd.tolist(), ch):
        bc.update((s, x + dx, z + dz) for dz in (-1, 0, 1) for dx in (-1, 0, 1))
        cols.update((s, 4 * x + k % 5, 4 * z + k // 5) for k in range(25))
        for e in range(27 * 19):
            qx, qz = 4 * x - 16 + e % 27, 4 * z - 8 + e // 27
            gk.add((s, qx, qz, 4 * x - 4 <= qx <= 4 * x + 6 and 4 * z - 4 <= qz <= 4 * z + 6))
        gk.update((s, 4 * x + 4 * (j & 1), 4 * z + 4 * (j >> 1), True) for j in range(4))
    u = gen_host.stage_units(sd, ch)
    assert u == {'chunks': len(ch), 'biome_chunks': len(bc), 'columns': len(cols), 'grid_keys': len(gk)}


def test_host_digest_folds_groups_like_the_fixture_writer():
    """mct.digest_host over uint64 group arrays (the stress test's folds of per-chunk digests)
    and over uint8 answers equals the fold tests/golden/make_digests.py wrote the C5 groups with."""
    import importlib.util
    import os
    import mct
    spec = importlib.util.spec_from_file_location('make_digests', os.path.join(os.path.dirname(__file__), 'golden', 'make_digests.py'))
    md = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(md)
    rng = np.random.default_rng(7)
    words = rng.integers(0, 2 ** 63, (6, 1024), dtype=np.uint64)
    answers = rng.integers(0, 256, (4, 4096), dtype=np.uint8)
    assert np.array_equal(mct.digest_host(words), md.fold(words))
    assert np.array_equal(mct.digest_host(answers), md.fold(answers))
