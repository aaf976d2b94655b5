"""Runs the reference's oracle-set checks (tests/test_chunks.cpp L2, tests/test_terrain.cpp)
against this library on the GPU: per seed of a `mccuda-oracle/1` set, the core configuration's
final density (bit for bit) and core fill against density/ and fill/, and the TERRAIN step's
biomes, aquifer fill and terrain against terrain/ when the set has them. Palettes are checked
against the compiled-in model first.

  python oracle_check.py DIR [--quick]       exit status 0 when every check passes
"""
import argparse
import json
import sys

import numpy as np

import mct
import oracle_data

CORE_PALETTE = ['minecraft:air', 'minecraft:stone', 'minecraft:water', 'minecraft:lava']


class Check:
    """Counts of one comparison, with the first mismatch kept for the report."""

    def __init__(self, name):
        self.name, self.total, self.failed, self.first = name, 0, 0, None

    def compare(self, want, got, where):
        want, got = np.asarray(want), np.asarray(got)
        bad = np.nonzero(want.reshape(-1) != got.reshape(-1))[0]
        self.total += want.size
        if len(bad):
            self.failed += len(bad)
            if self.first is None:
                i = int(bad[0])
                self.first = '%s: element %d want %s got %s' % (where, i, want.reshape(-1)[i], got.reshape(-1)[i])

    def ok(self):
        return self.failed == 0 and self.total > 0

    def line(self):
        return '%-28s %12d checked %10d mismatched%s' % (self.name, self.total, self.failed, ('  first: ' + self.first) if self.first else '')


def model_palette():
    tm = json.load(open(mct.TERRAIN_PATH))
    return list(tm['palette']), list(tm['biome_names'])


def run(directory, quick=False, batch=1024, log=print):
    s = oracle_data.load_oracle_set(directory)
    palette, biome_names = model_palette()
    if s.palette and s.palette != CORE_PALETTE:
        raise oracle_data.OracleError('core fill palette %r differs from %r' % (s.palette, CORE_PALETTE))
    if s.terrain_manifest is not None:
        if list(s.terrain_manifest['palette']) != palette:
            raise oracle_data.OracleError('terrain palette differs from the compiled-in model')
        if 'biome_names' in s.terrain_manifest and list(s.terrain_manifest['biome_names']) != biome_names:
            raise oracle_data.OracleError('biome names differ from the compiled-in model')
    bad_files = s.verify_hashes()
    if bad_files:
        log('files whose SHA-256 differs from the manifest: %s' % ', '.join(bad_files[:5]))
    checks = {k: Check(k) for k in ('density', 'core fill', 'biomes', 'aquifer fill', 'terrain')}
    g = mct.FullTerrainGenerator(max_batch_chunks=batch, pipeline_slots=2)
    for seed in s.seeds:
        if quick and not seed.quick:
            continue
        xz = seed.xz()
        d, f = g.generate_core(seed.seed, xz)
        checks['density'].compare(s.density(seed).view(np.uint32), d.view(np.uint32), 'seed#%d density' % seed.index)
        checks['core fill'].compare(s.fill(seed), f, 'seed#%d fill' % seed.index)
        if s.terrain_manifest is not None:
            wb, wa, wt = s.terrain(seed)
            g.set_seed(seed.seed)
            t, b, a = g.generate(xz, aquifer=True)
            checks['biomes'].compare(wb, b, 'seed#%d biomes' % seed.index)
            checks['aquifer fill'].compare(wa, a.astype(np.uint16), 'seed#%d aquifer' % seed.index)
            checks['terrain'].compare(wt, t.astype(np.uint16), 'seed#%d terrain' % seed.index)
    g.close()
    ran = [c for c in checks.values() if c.total > 0]
    for c in ran:
        log(c.line())
    return all(c.ok() for c in ran) and not bad_files, checks, bad_files


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument('dir')
    ap.add_argument('--quick', action='store_true', help='only the seeds the manifest marks quick')
    a = ap.parse_args()
    ok, _, _ = run(a.dir, quick=a.quick)
    sys.exit(0 if ok else 1)


if __name__ == '__main__':
    main()
